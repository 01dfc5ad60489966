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
