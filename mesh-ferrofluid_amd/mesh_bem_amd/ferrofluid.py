"""Callers of the BEM path in the mesh ferrofluid time step.

``MeshHelmholtzDecomposition`` mirrors the reference class
(``src/mesh_ferrofluid/mesh_Hemholtz_decomposition.py:18-143``): same
constructor, same ``solve(verts, faces, vert_vel) -> new_vel``.  The
reference builds [B, V, F, 1, Q^2, 3] tensors; here the dense sum over
(target vertex, face, quadrature point) is one HIP kernel
(``mba_helmholtz_decomp``) and only O(V) glue stays in torch:

    u = sum_f sum_q W_fq [ (n_f . u_fq) grad G + (n_f x u_fq) x grad G ]   (u1 + u2)
    new_vel = (I - n n^T) vel + (n . u) n                                  (vertex normals n)
"""

import torch

from . import _lib
from .geometry import compute_vert_normals
from .quadrature import get_gaussion_integration_points_and_weights


def tensor_rule(gaussQR: int, dtype=torch.float64, device="cuda") -> torch.Tensor:
    """[Q^2, 3] rows (r1, r2, w) of the reference's tensor rule on the unit
    panel (``mesh_Hemholtz_decomposition.py:33-48``): r1[a,b] = p[b],
    r2[a,b] = p[a] p[b], w[a,b] = w[a] w[b], flattened a * Q + b, with the
    float32-rounded Gauss-Legendre nodes."""
    p, w = get_gaussion_integration_points_and_weights(gaussQR, device="cpu", dtype=dtype)
    r1 = p[None, :].repeat(gaussQR, 1)
    r2 = p[:, None] * r1
    ww = w * w[:, None]
    return torch.stack([r1.reshape(-1), r2.reshape(-1), ww.reshape(-1)], dim=1).to(device).contiguous()


def decomposition_velocity(verts: torch.Tensor, faces: torch.Tensor, vert_vel: torch.Tensor, gaussQR: int,
                           targets: torch.Tensor = None) -> torch.Tensor:
    """u1 + u2 of the reference at ``targets`` ([Nt, 3]; default the
    vertices) for velocity fields ``vert_vel`` [B, V, 3] (or [V, 3]).
    Precision follows ``verts`` (float32 or float64).  Returns [B, Nt, 3]."""
    lib = _lib.load()
    dt = verts.dtype
    if dt not in (torch.float32, torch.float64):
        raise ValueError("decomposition_velocity: float32 or float64 vertices")
    V = verts.reshape(-1, 3).contiguous()
    Fc = faces.reshape(-1, 3).to(torch.int32).contiguous()
    vel = vert_vel.to(dt)
    if vel.dim() == 2:
        vel = vel[None]
    vel = vel.contiguous()
    if vel.shape[-2:] != V.shape:
        raise ValueError("decomposition_velocity: vert_vel must be [B, V, 3]")
    tg = V if targets is None else targets.reshape(-1, 3).to(dt).contiguous()
    rule = tensor_rule(gaussQR, dtype=dt, device=V.device)
    _lib.require_device(V, Fc, vel, tg, rule)
    B, nt = vel.shape[0], tg.shape[0]
    out = torch.empty((B, nt, 3), dtype=dt, device=V.device)
    if nt == 0 or B == 0:
        return out
    if Fc.shape[0] == 0:
        return out.zero_()
    prec = _lib.prec_of(dt)
    nq = rule.shape[0]
    wsb = lib.mba_helmholtz_decomp_workspace_bytes(prec, nt, Fc.shape[0], nq, B)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=V.device)
    st = lib.mba_helmholtz_decomp(prec, _lib.ptr(tg), nt, _lib.ptr(V), _lib.ptr(Fc), V.shape[0], Fc.shape[0],
                                  _lib.ptr(vel), B, _lib.ptr(rule), nq, _lib.ptr(out), _lib.ptr(ws), wsb,
                                  _lib.stream())
    _lib.check(st, "mba_helmholtz_decomp")
    return out


class MeshHelmholtzDecomposition(object):
    """Drop-in for ``src/mesh_ferrofluid/mesh_Hemholtz_decomposition.py:18-143``."""

    def __init__(self, gaussQR: int, device: torch.device = torch.device("cpu"), dtype: torch.dtype = torch.float32):
        self.gaussQR = gaussQR
        self.device = device
        self.dtype = dtype

    def solve(self, verts: torch.Tensor, faces: torch.Tensor, vert_vel: torch.Tensor) -> torch.Tensor:
        """verts [V, 3], faces [F, 3], vert_vel [B, V, 3] -> new_vel [B, V, 3]:
        tangential part of the velocity kept, normal part replaced by the
        normal component of the decomposition velocity (Da et al. 2016,
        explicit form; the reference's step 3)."""
        u = decomposition_velocity(verts, faces, vert_vel, self.gaussQR)  # [B, V, 3]
        vn = compute_vert_normals(verts, faces)  # [V, 3]
        un = (vn * u).sum(dim=-1, keepdim=True)
        vel = vert_vel.to(verts.dtype)
        return vel - (vel * vn).sum(dim=-1, keepdim=True) * vn + un * vn


# ---------------------------------------------------------------- pressure step
TRIPLE_JUNCTION = 6  # VertexType.TRIPLE_JUNCTION of src/utils/types.py:66-74


class MeshPressure:
    """Drop-in for ``src/mesh_ferrofluid/mesh_pressure.py:20-163``: same
    constructor and ``solve(verts, faces, dt, vel, vert_bc_flags, P_mag=None)
    -> new_vel``.

    Per step: vertex normals, pytorch3d's cot-Laplacian mean curvature and the
    Dirichlet / Neumann data (one kernel, ``mba_pressure_bc``), the Laplace
    BEM solve of ``MeshBEMSolver3d.solve`` (the reference's rule, Dirichlet
    vertices), then grad p = dp/dn n + the tangential gradient of p and
    new_vel = vel - (dt / rho) grad p (one kernel, ``mba_pressure_update``)."""

    def __init__(self, density_fluid: float, gravity_strength: float, surface_tension, contact_angle,
                 gaussQR: int, order_type: int, logger=None, device=torch.device("cpu"),
                 dtype: torch.dtype = torch.float32, *, bem_mode: int = None) -> None:
        """bem_mode (keyword-only, new): AssemblyMode of the Laplace solve;
        default the reference's rule (REF), NS for the NS P1 layers."""
        from .logger import Logger
        from .solver import MeshBEMSolver3d
        from .types import BEMType

        self.gaussQR = gaussQR
        self.order_type = order_type
        self.logger = logger if logger is not None else Logger(quiet=True)
        self.device = device
        self.dtype = dtype
        self.density_fluid = float(density_fluid)
        self.gravity_strength = float(gravity_strength)
        self.gravity = torch.tensor([0.0, -gravity_strength, 0.0], dtype=dtype, device=device)
        self.surface_tension = float(surface_tension)
        self.contact_angle = contact_angle
        from .types import AssemblyMode

        mode = int(AssemblyMode.REF) if bem_mode is None else int(bem_mode)
        self.BEM_solver = MeshBEMSolver3d(gaussQR=gaussQR, order_type=order_type, BEM_type=int(BEMType.LAPLACE),
                                          wavenumber=1.0, logger=self.logger, device=device, dtype=dtype,
                                          mode=mode)
        self.last = {}

    def boundary_data(self, verts: torch.Tensor, faces: torch.Tensor, dt: float, vel: torch.Tensor,
                      vert_bc_flags: torch.Tensor, P_mag: torch.Tensor = None):
        """(vertex normals [V, 3], kappa [V], Dirichlet p [B, V], Neumann dp/dn [B, V])."""
        from .mom import vertex_incidence

        lib = _lib.load()
        rdt = verts.dtype
        dev = verts.device
        V, F = verts.shape[0], faces.shape[0]
        B = vel.shape[0]
        ft = faces.to(torch.int32).contiguous()
        vptr, vidx = vertex_incidence(ft.cpu().numpy(), V)
        self._vinc = (torch.from_numpy(vptr).to(dev), torch.from_numpy(vidx).to(dev))
        u = vel.to(rdt).reshape(B, V, 3).contiguous()
        fl = vert_bc_flags.reshape(B, V).to(torch.uint8).contiguous()
        pm = None
        if P_mag is not None:
            pm = P_mag.to(rdt).reshape(-1, V).expand(B, V).contiguous()
        vn = torch.empty((V, 3), dtype=rdt, device=dev)
        kap = torch.empty(V, dtype=rdt, device=dev)
        pd = torch.empty((B, V), dtype=rdt, device=dev)
        dn = torch.empty((B, V), dtype=rdt, device=dev)
        st = lib.mba_pressure_bc(_lib.prec_of(rdt), _lib.ptr(verts.contiguous()), _lib.ptr(ft), V, F,
                                 _lib.ptr(self._vinc[0]), _lib.ptr(self._vinc[1]), _lib.ptr(u), _lib.ptr(fl),
                                 _lib.ptr(pm) if pm is not None else None, B, self.density_fluid,
                                 self.gravity_strength, self.surface_tension, float(dt), _lib.ptr(vn), _lib.ptr(kap),
                                 _lib.ptr(pd), _lib.ptr(dn), _lib.stream())
        _lib.check(st, "mba_pressure_bc")
        return vn, kap, pd, dn

    def solve(self, verts: torch.Tensor, faces: torch.Tensor, dt: float, vel: torch.Tensor,
              vert_bc_flags: torch.Tensor, P_mag: torch.Tensor = None) -> torch.Tensor:
        if (vert_bc_flags == TRIPLE_JUNCTION).sum().item() > 0:
            self.logger.ErrorLog("Not implemented for triple-junction points")
        lib = _lib.load()
        rdt = verts.dtype
        V, F = verts.shape[0], faces.shape[0]
        B = vel.shape[0]
        vn, kap, pd, dn = self.boundary_data(verts, faces, dt, vel, vert_bc_flags, P_mag)
        vert_dpdn, vert_p = self.BEM_solver.solve(verts=verts, faces=faces, vert_bc_Dirichlet=pd.reshape(B, V, 1),
                                                   vert_bc_Neumann=dn.reshape(B, V, 1),
                                                   vert_bc_flags=vert_bc_flags.reshape(B, V, 1), Amat_scaler=1.0)
        p = vert_p.reshape(B, V).to(rdt).contiguous()
        q = vert_dpdn.reshape(B, V).to(rdt).contiguous()
        u = vel.to(rdt).reshape(B, V, 3).contiguous()
        out = torch.empty_like(u)
        ft = faces.to(torch.int32).contiguous()
        st = lib.mba_pressure_update(_lib.prec_of(rdt), _lib.ptr(verts.contiguous()), _lib.ptr(ft), V, F,
                                     _lib.ptr(self._vinc[0]), _lib.ptr(self._vinc[1]), _lib.ptr(vn), _lib.ptr(p),
                                     _lib.ptr(q), _lib.ptr(u), B, self.density_fluid, float(dt), _lib.ptr(out),
                                     _lib.stream())
        _lib.check(st, "mba_pressure_update")
        self.last = {"kappa": kap, "p": p, "dpdn": q, "vert_normals": vn}
        return out.to(vel.dtype)
