"""2-D boundary elements on the HIP library: drop-in ``MeshBEMSolver2d``
(src/mesh_BEM/mesh_BEM_solver_2d.py:20-250).

The boundary is a closed curve of line segments: ``verts`` [V, 2], ``faces``
[F, 2] (int32 vertex pairs).  ``layer2d`` assembles the unpacked Galerkin
single (layer 1) or double (layer 2) layer [F, F, nphi, nphi] the way
``double_integration_on_panels_2d`` does (src/mesh_BEM/utils/integrations_2d.py:56-232),
``pack2d`` forms the [V, V] matrix (pack_matrix_tight, src/utils/BEM_utils.py:314-357),
``MeshBEMSolver2d.solve`` solves Dirichlet data for the Neumann trace:

  LAPLACE / HELMHOLTZ_BEM:  V t = 1/2 M g + K^T g
  HELMHOLTZ_MOM:            V t = M g
  (rhs_layer_2d.py:66-130, mesh_BEM_solver_2d.py:148-171)

with the dense solve on the LU kernels (complex64 for float32 input, complex128
for float64).  Neumann and mixed data raise NotImplementedError: the
reference's 2-D solver never builds the hypersingular layer it would need
(mesh_BEM_solver_2d.py:43-75 constructs single, double and rhs layers only)
and its right-hand side reads the double layer only inside the Dirichlet
branch (rhs_layer_2d.py:106-147), so those cases fail there too.
"""

from typing import List, Optional

import torch

from . import _lib
from .geometry import compute_vert_areas
from .logger import Logger
from .mom import vertex_incidence
from .quadrature import get_gaussion_integration_points_and_weights
from .types import BEMType, BoundaryType, OrderType


def _check_curve(verts: torch.Tensor, faces: torch.Tensor):
    if verts.dim() != 2 or verts.shape[1] != 2:
        raise ValueError("2-D BEM needs verts [V, 2]")
    if faces.dim() != 2 or faces.shape[1] != 2:
        raise ValueError("2-D BEM needs segment faces [F, 2]")
    if verts.device.type != "cuda":
        raise ValueError("2-D BEM runs on the GPU: verts must be a device tensor")


def layer2d(verts: torch.Tensor, faces: torch.Tensor, gaussQR: int, layer: int, order_type: int,
            BEM_type: int = int(BEMType.LAPLACE), wavenumber: float = 1.0) -> torch.Tensor:
    """Unpacked [F, F, nphi, nphi] single (layer 1) or double (layer 2) layer;
    real for LAPLACE, complex for the Helmholtz types; dtype of verts."""
    _check_curve(verts, faces)
    lib = _lib.load()
    dt = verts.dtype
    prec = _lib.prec_of(dt)
    dev = verts.device
    gp, gw = get_gaussion_integration_points_and_weights(gaussQR, dtype=dt)
    gp, gw = gp.to(dev).contiguous(), gw.to(dev).contiguous()
    F = faces.shape[0]
    nphi = 2 if order_type == int(OrderType.LINEAR) else 1
    cplx = BEMType.use_complex(BEM_type)
    odt = (torch.complex64 if dt == torch.float32 else torch.complex128) if cplx else dt
    out = torch.empty((F, F, nphi, nphi), dtype=odt, device=dev)
    ft = faces.to(torch.int32).contiguous()
    st = lib.mba_bem2d_layer(prec, int(BEM_type), int(layer), int(order_type), _lib.ptr(verts.contiguous()),
                             _lib.ptr(ft), verts.shape[0], F, _lib.ptr(gp), _lib.ptr(gw), gp.numel(),
                             float(wavenumber), _lib.ptr(out), _lib.stream())
    _lib.check(st, "mba_bem2d_layer")
    return out


def pack2d(unpacked: torch.Tensor, faces: torch.Tensor, n_verts: int) -> torch.Tensor:
    """[F, F, nphi, nphi] -> [V, V] (fixed-order gather over the vertex slots)."""
    lib = _lib.load()
    F = faces.shape[0]
    nphi = unpacked.shape[-1]
    cplx = unpacked.is_complex()
    prec = _lib.prec_of(unpacked.real.dtype if cplx else unpacked.dtype)
    vptr, vidx = vertex_incidence(faces.detach().cpu().numpy(), n_verts)
    dev = unpacked.device
    vp = torch.from_numpy(vptr).to(dev)
    vi = torch.from_numpy(vidx).to(dev)
    out = torch.empty((n_verts, n_verts), dtype=unpacked.dtype, device=dev)
    st = lib.mba_bem2d_pack(prec, int(cplx), _lib.ptr(unpacked.contiguous()), F, nphi, _lib.ptr(vp), _lib.ptr(vi),
                            n_verts, _lib.ptr(out), _lib.stream())
    _lib.check(st, "mba_bem2d_pack")
    return out


def dirichlet_rhs2d(K_tight: Optional[torch.Tensor], g: torch.Tensor, vert_areas: torch.Tensor, c_mass: float,
                    c_k: float, dtype: torch.dtype) -> torch.Tensor:
    """rhs = c_mass g * va + c_k K^T g (dtype: the matrices' dtype)."""
    lib = _lib.load()
    cplx = dtype.is_complex
    rdt = torch.float32 if dtype in (torch.float32, torch.complex64) else torch.float64
    V = vert_areas.numel()
    gv = g.reshape(-1).to(dtype).contiguous()
    va = vert_areas.reshape(-1).to(rdt).contiguous()
    out = torch.empty(V, dtype=dtype, device=va.device)
    kp = _lib.ptr(K_tight.contiguous()) if K_tight is not None else None
    st = lib.mba_bem2d_dirichlet_rhs(_lib.prec_of(rdt), int(cplx), kp, V, _lib.ptr(gv), _lib.ptr(va),
                                     float(c_mass), float(c_k), _lib.ptr(out), _lib.stream())
    _lib.check(st, "mba_bem2d_dirichlet_rhs")
    return out


def _dense_solve(A: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """A x = b on the LU kernels (a real system goes through the complex kernels
    with zero imaginary parts)."""
    from .dense import LUFactor

    real = not A.is_complex()
    cdt = (torch.complex64 if A.dtype == torch.float32 else torch.complex128) if real else A.dtype
    x = LUFactor(A.to(cdt), overwrite=True).solve(b.to(cdt))
    return x.real.contiguous() if real else x


class MeshBEMSolver2d:
    """Drop-in for src/mesh_BEM/mesh_BEM_solver_2d.py:20-250 (same constructor
    and ``solve`` signature and return values)."""

    rank = 2

    def __init__(
        self,
        gaussQR: int,
        order_type: int,
        BEM_type: int,
        wavenumber: float,
        logger: Logger,
        dtype: torch.dtype = torch.float32,
        device: torch.device = torch.device("cpu"),
    ) -> None:
        self.gaussQR = gaussQR
        self.order_type = order_type
        self.BEM_type = BEM_type
        self.wavenumber = wavenumber
        self.logger = logger if logger is not None else Logger(quiet=True)
        self.dtype = dtype
        self.device = device
        # raises for orders the reference's table does not hold
        self.gaussian_points_1d, self.gaussian_weights_1d = get_gaussion_integration_points_and_weights(
            gaussQR, device=device, dtype=dtype)

    def solve(
        self,
        verts: torch.Tensor,
        faces: torch.Tensor,
        vert_bc_Dirichlet: torch.Tensor,
        vert_bc_Neumann: torch.Tensor,
        vert_bc_flags: torch.Tensor,
        Amat_scaler: float = None,
    ) -> List[torch.Tensor]:
        """Dirichlet data -> (Neumann [B, V, 1], Dirichlet [B, V, 1])."""
        _check_curve(verts, faces)
        V = verts.shape[-2]
        B = vert_bc_flags.shape[0]
        faces = faces.to(torch.int32)
        va = compute_vert_areas(verts, faces, keepdim=False)
        mom = self.BEM_type == int(BEMType.HELMHOLTZ_MOM)
        Vt = Kt = None
        neu, dir_ = [], []
        for b in range(B):
            flags = vert_bc_flags[b].reshape(-1)
            nD = int((flags == int(BoundaryType.DIRICHLET)).sum())
            nN = int((flags == int(BoundaryType.NEUMANN)).sum())
            self.logger.InfoLog(f"Batch [{b}]: (Vertex attached) number of Dirichlet B.C. = {nD}, "
                                f"number of Neumann B.C. = {nN}, number of whole zone = {flags.numel()}")
            if nN > 0:
                raise NotImplementedError(
                    "2-D Neumann / mixed data need the hypersingular layer, which the 2-D solver does not build")
            if nD != V:
                raise ValueError(f"batch {b}: every vertex flag must be DIRICHLET or NEUMANN")
            if Vt is None:
                Vt = pack2d(layer2d(verts, faces, self.gaussQR, 1, self.order_type, self.BEM_type,
                                    self.wavenumber), faces, V)
                if not mom:
                    Kt = pack2d(layer2d(verts, faces, self.gaussQR, 2, self.order_type, self.BEM_type,
                                        self.wavenumber), faces, V)
                if Amat_scaler is not None:
                    Vt = Amat_scaler * Vt
            g = vert_bc_Dirichlet[b].reshape(-1)
            rhs = dirichlet_rhs2d(Kt, g, va, 1.0 if mom else 0.5, 0.0 if mom else 1.0, Vt.dtype)
            t = _dense_solve(Vt, rhs)
            neu.append(t.reshape(V, 1))
            dir_.append(vert_bc_Dirichlet[b].reshape(V, 1))
        return torch.stack(neu, 0), torch.stack(dir_, 0)


def ellipse_curve(n: int, a: float = 1.0, b: float = 1.0):
    """Closed counter-clockwise polygon of n segments on the ellipse
    x^2/a^2 + y^2/b^2 = 1 (the reference's 2-D demo builds the circle the same
    way, demo/demo_Laplace_2d.py:100-110): (verts [n, 2] float64 numpy,
    faces [n, 2] int32 numpy)."""
    import numpy as np

    th = 2.0 * np.pi * np.arange(n) / n
    verts = np.stack([a * np.cos(th), b * np.sin(th)], axis=1)
    faces = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1).astype(np.int32)
    return verts, faces
