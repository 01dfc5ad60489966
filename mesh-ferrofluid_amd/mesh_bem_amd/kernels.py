"""Op-level drop-ins for the reference's kernel wrappers.

Same class names, ``forward`` signatures, output shapes and dtypes as
src/mesh_BEM/kernel_functions.py:26-171: ``MeshMOMZmatKernel.apply`` returns
the complex ``[1, 3F, 3F]`` face-edge matrix, ``MeshMOMRhsKernel.apply`` the
complex ``[1, F, 3, 1]`` RHS and ``MeshMOMInterpolateJsKernel.apply`` the
complex ``[B, 3F, 3]`` face-vertex currents.  As in the reference there is no
backward.  The work runs in the HIP library (include/mesh_bem_amd.h) on
torch's current stream; there is no CPU path.
"""

from typing import Any

import torch

from . import _lib


def _cdtype(dtype):
    return torch.complex64 if dtype == torch.float32 else torch.complex128


def _check_mesh(verts, faces, face_areas, face_normals=None):
    """Operand shapes the kernels index by: verts [V, 3], faces [F, 3], one
    area (and one normal) per face.  Raises ValueError before any launch."""
    if verts.ndim != 2 or verts.shape[1] != 3 or faces.ndim != 2 or faces.shape[1] != 3:
        raise ValueError("verts must be [V, 3] and faces [F, 3]")
    F = faces.shape[0]
    if face_areas.numel() != F:
        raise ValueError(f"face_areas must hold one value per face ({F}), got {face_areas.numel()}")
    if face_normals is not None and face_normals.numel() != 3 * F:
        raise ValueError(f"face_normals must be [F, 3] (F = {F}), got {tuple(face_normals.shape)}")


def _check_rule(points, weights):
    if points.numel() != weights.numel() or points.numel() == 0:
        raise ValueError("gaussian points and weights must be non-empty and of equal length")


class MeshMOMZmatKernel(torch.autograd.Function):
    @staticmethod
    def forward(
        ctx,
        verts: torch.Tensor,
        faces: torch.Tensor,
        face_areas: torch.Tensor,
        face_normals: torch.Tensor,
        gaussian_points_1d_x: torch.Tensor,
        gaussian_weights_1d_x: torch.Tensor,
        gaussian_points_1d_y: torch.Tensor,
        gaussian_weights_1d_y: torch.Tensor,
        wavenumber: float,
    ) -> Any:
        _check_mesh(verts, faces, face_areas, face_normals)
        _check_rule(gaussian_points_1d_x, gaussian_weights_1d_x)
        _check_rule(gaussian_points_1d_y, gaussian_weights_1d_y)
        lib = _lib.load()
        N_faces, dim = faces.shape
        dt = verts.dtype
        args = [verts, faces.to(torch.int32), face_areas.to(dt), face_normals.to(dt),
                gaussian_points_1d_x.to(dt), gaussian_weights_1d_x.to(dt),
                gaussian_points_1d_y.to(dt), gaussian_weights_1d_y.to(dt)]
        args = [a.contiguous() for a in args]
        _lib.require_device(*args)
        out = torch.empty((1, N_faces * dim, N_faces * dim), dtype=_cdtype(dt), device=verts.device)
        st = lib.mba_mom_zmat_faceedge_ref(
            _lib.prec_of(dt), *[_lib.ptr(a) for a in args[:6]], args[4].numel(),
            _lib.ptr(args[6]), _lib.ptr(args[7]), args[6].numel(), float(wavenumber),
            verts.shape[0], N_faces, _lib.ptr(out), 3 * N_faces, _lib.stream())
        _lib.check(st, "mba_mom_zmat_faceedge_ref")
        return out


class MeshMOMRhsKernel(torch.autograd.Function):
    @staticmethod
    def forward(
        ctx,
        verts: torch.Tensor,
        faces: torch.Tensor,
        face_areas: torch.Tensor,
        face_normals: torch.Tensor,
        gaussian_points_1d: torch.Tensor,
        gaussian_weights_1d: torch.Tensor,
        wavenumber: float,
    ) -> Any:
        _check_mesh(verts, faces, face_areas, face_normals)
        _check_rule(gaussian_points_1d, gaussian_weights_1d)
        lib = _lib.load()
        N_faces, dim = faces.shape
        dt = verts.dtype
        args = [verts, faces.to(torch.int32), face_areas.to(dt), face_normals.to(dt),
                gaussian_points_1d.to(dt), gaussian_weights_1d.to(dt)]
        args = [a.contiguous() for a in args]
        _lib.require_device(*args)
        out = torch.empty((1, N_faces, dim, 1), dtype=_cdtype(dt), device=verts.device)
        st = lib.mba_mom_rhs_faceedge_ref(
            _lib.prec_of(dt), *[_lib.ptr(a) for a in args], args[4].numel(), float(wavenumber),
            verts.shape[0], N_faces, _lib.ptr(out), _lib.stream())
        _lib.check(st, "mba_mom_rhs_faceedge_ref")
        return out


class MeshMOMInterpolateJsKernel(torch.autograd.Function):
    @staticmethod
    def forward(
        ctx,
        verts: torch.Tensor,
        faces: torch.Tensor,
        face_edge_indices: torch.Tensor,
        face_areas: torch.Tensor,
        edge_I_coeff: torch.Tensor,
    ) -> Any:
        _check_mesh(verts, faces, face_areas)
        if face_edge_indices.numel() != 3 * faces.shape[0]:
            raise ValueError("face_edge_indices must hold 3 entries per face")
        lib = _lib.load()
        N_faces, dim = faces.shape
        dt = verts.dtype
        B = edge_I_coeff.shape[0]
        icoef = edge_I_coeff.reshape(B, -1).to(_cdtype(dt)).contiguous()
        args = [verts.contiguous(), faces.to(torch.int32).contiguous(),
                face_edge_indices.to(torch.int32).contiguous(), face_areas.to(dt).contiguous()]
        _lib.require_device(*args, icoef)
        out = torch.empty((B, N_faces * dim, dim), dtype=_cdtype(dt), device=verts.device)
        st = lib.mba_mom_interp_faceedge(
            _lib.prec_of(dt), *[_lib.ptr(a) for a in args], _lib.ptr(icoef), B, icoef.shape[1],
            verts.shape[0], N_faces, _lib.ptr(out), _lib.stream())
        _lib.check(st, "mba_mom_interp_faceedge")
        return out
