"""Drop-in ``MeshBEMSolver3d`` (src/mesh_BEM/mesh_BEM_solver_3d.py:31-373).

Constructor and method signatures are the reference's; new options are
keyword-only with defaults that reproduce the reference:

* ``mode``: ``AssemblyMode.REF`` (default) runs the reference rule (collapsed
  Gauss of order gaussQR / gaussQR-1, CFIE, fixed incidence);
  ``AssemblyMode.NS`` runs the EFIE with the 7-point rule, analytic FP64
  near-singular terms and the plane waves given by ``incidence``.
* ``precision``: ``None`` solves in the dtype of the vertices (as the reference:
  complex64 for float32 input); ``"fp64"`` assembles and factors in FP64;
  ``"mixed"`` assembles the FP64 matrix, factors its FP32 rounding on MFMA and
  refines with FP64 residuals to ``refine_tol``.
* ``incidence``: (directions [D,3], polarisations [D,3]) for NS mode; default
  x-polarised along +z.

``solve_MOM`` returns the vertex currents [1, V, 3] (complex) like the
reference; ``solve_MOM_rwg`` returns the RWG coefficients and a record of the
steps.  ``solve`` is the Laplace Dirichlet solve (src/mesh_BEM/mesh_BEM_solver_3d.py:197-373).
"""

import time
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .dense import LUFactor, refine_solve, to_c64
from .geometry import compute_face_areas, compute_face_normals
from .kernels import MeshMOMRhsKernel, MeshMOMZmatKernel
from .logger import Logger
from .mom import MomMesh, assemble_ns, currents_to_vertices, rhs_ns
from .quadrature import get_gaussion_integration_points_and_weights
from .types import AssemblyMode, BEMType


class MeshBEMSolver3d:
    rank = 3

    def __init__(
        self,
        gaussQR: int,
        order_type: int,
        BEM_type: int,
        wavenumber: float,
        logger: Logger,
        dtype: torch.dtype = torch.float32,
        device: torch.device = torch.device("cpu"),
        *,
        mode: int = int(AssemblyMode.REF),
        precision: Optional[str] = None,
        incidence=None,
        refine_tol: float = 1e-8,
        verbose: bool = False,
    ) -> None:
        self._gaussQR = gaussQR
        self._order_type = order_type
        self._BEM_type = BEM_type
        self._wavenumber = wavenumber
        self.logger = logger if logger is not None else Logger(quiet=True)
        self.dtype = dtype
        self.device = device
        self.mode = int(mode)
        if precision not in (None, "fp32", "fp64", "mixed"):
            raise ValueError(f"precision must be None, 'fp32', 'fp64' or 'mixed', not {precision!r}")
        self.precision = precision
        self.incidence = incidence
        self.refine_tol = refine_tol
        self.verbose = verbose
        self.last_record = {}
        self.gaussian_points_1d_x, self.gaussian_weights_1d_x = get_gaussion_integration_points_and_weights(
            gaussQR, device=device, dtype=dtype)
        if gaussQR - 1 >= 1:
            self.gaussian_points_1d_y, self.gaussian_weights_1d_y = \
                get_gaussion_integration_points_and_weights(gaussQR - 1, device=device, dtype=dtype)
        # RHS table of the reference (src/mesh_BEM/rhs_layer/rhs_layer_3d.py:49-53)
        self.rhs_gaussian_points_1d = self.gaussian_points_1d_x
        self.rhs_gaussian_weights_1d = self.gaussian_weights_1d_x
        self._laplace = None

    # -------------------------------------------------------------- properties
    @property
    def gaussQR(self) -> int:
        return self._gaussQR

    @property
    def order_type(self) -> int:
        return self._order_type

    @property
    def BEM_type(self) -> int:
        return self._BEM_type

    @property
    def wavenumber(self) -> float:
        return self._wavenumber

    def InfoLog(self, *args, **kwargs):
        return self.logger.InfoLog(*args, **kwargs)

    def WarnLog(self, *args, **kwargs):
        return self.logger.WarnLog(*args, **kwargs)

    def ErrorLog(self, *args, **kwargs):
        return self.logger.ErrorLog(*args, **kwargs)

    # -------------------------------------------------------------- MoM
    def _assembly_dtype(self, verts_dtype):
        if self.precision in ("fp64", "mixed"):
            return torch.float64
        if self.precision == "fp32":
            return torch.float32
        return verts_dtype

    def _system(self, mesh: MomMesh, vt: torch.Tensor, faces: torch.Tensor):
        """(Z [E,E], V [E, D]) in the assembly precision."""
        k = self.wavenumber
        E = mesh.n_edges
        prec = _lib.prec_of(vt.dtype)
        if self.mode == int(AssemblyMode.REF):
            fa = compute_face_areas(vt, faces, keepdim=True)
            fn = compute_face_normals(vt, faces)
            gpx = self.gaussian_points_1d_x.to(vt.dtype)
            gwx = self.gaussian_weights_1d_x.to(vt.dtype)
            gpy = self.gaussian_points_1d_y.to(vt.dtype)
            gwy = self.gaussian_weights_1d_y.to(vt.dtype)
            zf = MeshMOMZmatKernel.apply(vt, faces, fa, fn, gpx, gwx, gpy, gwy, k)[0]
            Z = torch.empty((E, E), dtype=zf.dtype, device=zf.device)
            lib = _lib.load()
            st = lib.mba_mom_gather_rwg(prec, _lib.ptr(zf), zf.stride(0), 0, zf.shape[1],
                                        _lib.ptr(mesh.rwg_fe), E, _lib.ptr(Z), Z.stride(0),
                                        _lib.stream())
            _lib.check(st, "mba_mom_gather_rwg")
            vf = MeshMOMRhsKernel.apply(vt, faces, fa, fn, gpx, gwx, k).reshape(-1, 1).contiguous()
            V = torch.empty((E, 1), dtype=vf.dtype, device=vf.device)
            st = lib.mba_mom_gather_rhs(prec, _lib.ptr(vf), _lib.ptr(mesh.rwg_fe), E, 1, _lib.ptr(V),
                                        _lib.stream())
            _lib.check(st, "mba_mom_gather_rhs")
            return Z, V
        if self.incidence is None:
            dirs, pols = np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]])
        else:
            dirs, pols = self.incidence
        m = MomMesh(**{**mesh.__dict__, "verts": vt})
        Z = assemble_ns(m, k, prec)
        V = rhs_ns(m, k, prec, dirs, pols)
        return Z, V

    def solve_MOM_rwg(self, verts: torch.Tensor, faces: torch.Tensor, edges: torch.Tensor,
                      face_edge_indices: torch.Tensor, mesh: Optional[MomMesh] = None):
        """RWG coefficients I [E, D] (complex128 for fp64/mixed) and the record of the solve."""
        if self._BEM_type != int(BEMType.HELMHOLTZ_MOM):
            raise AssertionError("solve_MOM needs BEM_type == HELMHOLTZ_MOM")
        rec = {}
        sync = torch.cuda.synchronize
        t0 = time.perf_counter()
        if mesh is None:
            mesh = MomMesh.build(verts, faces, face_edge_indices, int(edges.shape[0]),
                                 device=verts.device, near=self.mode == int(AssemblyMode.NS))
        adt = self._assembly_dtype(verts.dtype)
        vt = mesh.verts.to(adt).contiguous()
        ft = mesh.faces
        Z, V = self._system(mesh, vt, ft)
        sync()
        t1 = time.perf_counter()
        rec["assembly_s"] = t1 - t0
        if self.precision == "mixed":
            lu = LUFactor(to_c64(Z), overwrite=True)
            sync()
            t2 = time.perf_counter()
            rec["factor_s"] = t2 - t1
            I, info = refine_solve(Z, V, lu32=lu, tol=self.refine_tol)
            sync()
            rec["refine_s"] = time.perf_counter() - t2
            rec["refine_iterations"] = info.iterations
            rec["refine_converged"] = info.converged
            rec["refine_fallback"] = info.fallback
        else:
            lu = LUFactor(Z, overwrite=True)
            sync()
            t2 = time.perf_counter()
            rec["factor_s"] = t2 - t1
            I = lu.solve(V)
            sync()
            rec["refine_s"] = time.perf_counter() - t2
        rec["solve_s"] = time.perf_counter() - t0
        self.last_record = rec
        return I, rec, mesh

    def solve_MOM(self, verts: torch.Tensor, faces: torch.Tensor, edges: torch.Tensor,
                  face_edge_indices: torch.Tensor) -> torch.Tensor:
        """Vertex currents [B, V, 3] (B = number of incident fields; 1 in REF mode)."""
        I, rec, mesh = self.solve_MOM_rwg(verts, faces, edges, face_edge_indices)
        icoef = I.T.contiguous().to(torch.complex128 if verts.dtype == torch.float64 else torch.complex64)
        m = MomMesh(**{**mesh.__dict__, "verts": verts.contiguous()})
        _, vj = currents_to_vertices(m, icoef)
        if self.verbose:
            self.InfoLog(f"solve_MOM: {rec}")
        return vj

    # -------------------------------------------------------------- Laplace
    def solve(self, verts: torch.Tensor, faces: torch.Tensor, vert_bc_Dirichlet: torch.Tensor,
              vert_bc_Neumann: torch.Tensor, vert_bc_flags: torch.Tensor,
              Amat_scaler: float = None) -> List[torch.Tensor]:
        from .laplace import laplace_solve

        return laplace_solve(self, verts, faces, vert_bc_Dirichlet, vert_bc_Neumann, vert_bc_flags,
                             Amat_scaler)
