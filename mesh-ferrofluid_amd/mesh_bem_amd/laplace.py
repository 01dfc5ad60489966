"""Laplace BEM on the HIP library.

Reference rule (the drop-in ``MeshBEMSolver3d.solve``,
src/mesh_BEM/mesh_BEM_solver_3d.py:197-373):

* ``relation_tables`` builds the point tables of the four panel relations the
  way the reference does: the rank-3 rule of
  src/mesh_BEM/single_layer/abstract_single_layer.py:49-63 (r1 = g_b,
  r2 = g_a g_b, weight w_a w_b) is tensorised into (xsi, eta1, eta2, eta3)
  (src/mesh_BEM/utils/integrations_3d.py:354-364) and mapped per relation
  (get_rands_and_jacobian_3d, integrations_3d.py:18-249).  The maps are
  applied to that input exactly as the reference applies them, including its
  second collapse of (xsi, eta1) and its three (of six) same-face regions.
* ``layer_ref`` assembles the unpacked [F, F, nphi, nphi] single or double
  layer on the GPU; ``pack`` forms the [V, V] matrix, ``dirichlet_rhs`` the
  right-hand side 1/2 M g + K^T g.

NS P1 rule (``MeshBEMSolver3d(..., mode=AssemblyMode.NS).solve``):
``p1_operators`` returns the unpacked P1 Galerkin single and double layer
(7-point rule, FP64 closed-form inner integrals on near pairs) in the layout
of ``layer_ref``; the Dirichlet system uses the consistent P1 mass matrix
(``mass_apply``) where the reference lumps it (rhs_layer_3d.py:127): with
the lumped mass the exact layers miss the demo's known answer by tens of
per cent, with the consistent one they converge to it.

NS rule for the magnetic scalar potential: ``p0_operators`` returns
the P0 Galerkin single layer V and adjoint double layer K' (7-point rule, FP64
analytic inner integrals on near pairs); ``magnetic_charge`` solves
[(1/chi + 1/2) M + K'] sigma = M (H0 . n) for the surface charge of a linear
magnetisable body in a uniform field.
"""

from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .geometry import compute_face_areas, compute_face_normals, compute_vert_areas
from .mom import near_pairs_gpu, rule_tensor, vertex_incidence
from .quadrature import get_gaussion_integration_points_and_weights
from .types import BEMType, BoundaryType, OrderType

# Same-face maps: rows of the three 4x4 matrices acting on
# (xsi, xsi eta1, xsi eta1 eta2, xsi eta1 eta2 eta3) (integrations_3d.py:204-229).
_SAME_FACE = np.array(
    [
        [[1, 0, 0, 0], [1, -1, 1, 0], [0, 0, 0, 1], [0, 0, 1, 0]],
        [[1, 0, 0, 0], [0, 1, -1, 1], [0, 0, 1, 0], [0, 0, 0, 1]],
        [[1, 0, 0, -1], [0, 1, 0, -1], [0, 0, 0, -1], [0, 0, 1, -1]],
    ],
    dtype=np.float64,
)


def _edge_maps(x, e1, e2, e3):
    """The five common-edge sub-domains as (w0, w1, w2, w3, jacobian) -> points
    (r1x, r2x, r1y, r2y) = (w0, w3, w0 + w1, w2)."""
    base = x ** 3 * e1 ** 2
    subs = [
        ((x, -x * e1 * e2, x * e1 * (1 - e2), x * e1 * e3), base),
        ((x, -x * e1 * e2 * e3, x * e1 * e2 * (1 - e3), x * e1), base * e2),
        ((x * (1 - e1 * e2), x * e1 * e2, x * e1 * e2 * e3, x * e1 * (1 - e2)), base * e2),
        ((x * (1 - e1 * e2 * e3), x * e1 * e2 * e3, x * e1, x * e1 * e2 * (1 - e3)), base * e2),
        ((x * (1 - e1 * e2 * e3), x * e1 * e2 * e3, x * e1 * e2, x * e1 * (1 - e2 * e3)), base * e2),
    ]
    return [((w0, w3, w0 + w1, w2), j) for (w0, w1, w2, w3), j in subs]


def relation_tables(gp: np.ndarray, gw: np.ndarray):
    """[table (rows, 5), offsets (5,)] for relations SEPARATE, SAME_VERTEX,
    SAME_EDGE, SAME_FACE; a row is (r1x, r2x, r1y, r2y, weight * jacobian)."""
    gp = np.asarray(gp, dtype=np.float64)
    gw = np.asarray(gw, dtype=np.float64)
    Q = gp.size
    # rank-3 rule, index a*Q + b
    r1 = np.broadcast_to(gp[None, :], (Q, Q)).reshape(-1)
    r2 = np.outer(gp, gp).reshape(-1)
    w2 = np.outer(gw, gw).reshape(-1)
    # tensorisation: index u*Q2 + v; (xsi, eta1) from v, (eta2, eta3) from u
    xsi = np.tile(r1, r1.size)
    eta1 = np.tile(r2, r2.size)
    eta2 = np.repeat(r1, r1.size)
    eta3 = np.repeat(r2, r2.size)
    wyx = np.repeat(w2, w2.size) * np.tile(w2, w2.size)
    groups = []
    groups.append([((xsi, eta1 * xsi, eta2, eta2 * eta3), xsi * eta2)])
    jv = xsi ** 3 * eta2
    groups.append([((xsi, xsi * eta1, xsi * eta2, xsi * eta2 * eta3), jv),
                   ((xsi * eta2, xsi * eta2 * eta3, xsi, xsi * eta1), jv)])
    groups.append(_edge_maps(xsi, eta1, eta2, eta3))
    w = np.stack([xsi, xsi * eta1, xsi * eta1 * eta2, xsi * eta1 * eta2 * eta3])
    jf = xsi ** 3 * eta1 ** 2 * eta2
    face = []
    for M in _SAME_FACE:
        z = M @ w
        face.append(((z[0], z[1], z[0] - z[2], z[1] - z[3]), jf))
    groups.append(face)
    rows, offsets = [], [0]
    for g in groups:
        blk = np.concatenate([np.stack([*pts, wyx * jac], axis=1) for pts, jac in g], axis=0)
        rows.append(blk)
        offsets.append(offsets[-1] + blk.shape[0])
    return np.ascontiguousarray(np.concatenate(rows, 0)), np.array(offsets, dtype=np.int64)


def _check_vf(verts, faces):
    if verts.ndim != 2 or verts.shape[1] != 3 or faces.ndim != 2 or faces.shape[1] != 3:
        raise ValueError("verts must be [V, 3] and faces [F, 3]")


def _rdt(prec):
    return torch.float32 if prec == _lib.MBA_F32 else torch.float64


def layer_ref(verts: torch.Tensor, faces: torch.Tensor, gaussQR: int, layer: int, order_type: int,
              tables=None) -> torch.Tensor:
    """Unpacked [F, F, nphi, nphi] Laplace layer (1 single, 2 double), dtype of verts."""
    _check_vf(verts, faces)
    lib = _lib.load()
    dt = verts.dtype
    prec = _lib.prec_of(dt)
    dev = verts.device
    F = faces.shape[0]
    if tables is None:
        gp, gw = get_gaussion_integration_points_and_weights(gaussQR, dtype=torch.float64)
        tables = relation_tables(gp.numpy(), gw.numpy())
    tab, off = tables
    tab_t = torch.from_numpy(tab).to(dt).to(dev)
    off_t = torch.from_numpy(off).to(dev)
    nphi = 3 if order_type == int(OrderType.LINEAR) else 1
    fa = compute_face_areas(verts, faces, keepdim=False).contiguous()
    fn = compute_face_normals(verts, faces).contiguous()
    ft = faces.to(torch.int32).contiguous()
    out = torch.empty((F, F, nphi, nphi), dtype=dt, device=dev)
    st = lib.mba_laplace_layer_ref(prec, _lib.ptr(verts.contiguous()), _lib.ptr(ft), _lib.ptr(fa),
                                   _lib.ptr(fn), verts.shape[0], F, _lib.ptr(tab_t), _lib.ptr(off_t),
                                   int(layer), nphi, _lib.ptr(out), _lib.stream())
    _lib.check(st, "mba_laplace_layer_ref")
    return out


def pack(unpacked: torch.Tensor, faces: torch.Tensor, n_verts: int) -> torch.Tensor:
    """[F, F, nphi, nphi] -> [V, V] by the vertex incidence (fixed summation order)."""
    F = faces.shape[0]
    nphi = unpacked.shape[-1]
    if unpacked.shape != (F, F, nphi, nphi) or nphi not in (1, 3):
        raise ValueError(f"unpacked must be [F, F, nphi, nphi] with F = {F}, nphi 1 or 3")
    lib = _lib.load()
    vptr, vidx = vertex_incidence(faces.detach().cpu().numpy(), n_verts)
    dev = unpacked.device
    vp = torch.from_numpy(vptr).to(dev)
    vi = torch.from_numpy(vidx).to(dev)
    out = torch.empty((n_verts, n_verts), dtype=unpacked.dtype, device=dev)
    st = lib.mba_laplace_pack(_lib.prec_of(unpacked.dtype), _lib.ptr(unpacked.contiguous()), F, nphi,
                              _lib.ptr(vp), _lib.ptr(vi), n_verts, _lib.ptr(out), _lib.stream())
    _lib.check(st, "mba_laplace_pack")
    return out


def dirichlet_rhs(K_tight: torch.Tensor, g: torch.Tensor, vert_areas: torch.Tensor) -> torch.Tensor:
    V = K_tight.shape[0]
    if K_tight.shape != (V, V) or g.numel() != V or vert_areas.numel() != V:
        raise ValueError(f"K must be [V, V] and g, vert_areas hold V = {V} values")
    lib = _lib.load()
    out = torch.empty(V, dtype=K_tight.dtype, device=K_tight.device)
    st = lib.mba_laplace_dirichlet_rhs(_lib.prec_of(K_tight.dtype), _lib.ptr(K_tight.contiguous()), V,
                                       _lib.ptr(g.reshape(-1).to(K_tight.dtype).contiguous()),
                                       _lib.ptr(vert_areas.reshape(-1).to(K_tight.dtype).contiguous()),
                                       _lib.ptr(out), _lib.stream())
    _lib.check(st, "mba_laplace_dirichlet_rhs")
    return out


def _real_solve(A: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Real dense solve through the complex LU kernels (imaginary parts zero)."""
    from .dense import LUFactor

    cdt = torch.complex64 if A.dtype == torch.float32 else torch.complex128
    f = LUFactor(A.to(cdt), overwrite=True)
    return f.solve(b.to(cdt)).real.contiguous()


def laplace_solve(solver, verts: torch.Tensor, faces: torch.Tensor, vert_bc_Dirichlet: torch.Tensor,
                  vert_bc_Neumann: torch.Tensor, vert_bc_flags: torch.Tensor,
                  Amat_scaler: Optional[float] = None) -> List[torch.Tensor]:
    """Dirichlet-to-Neumann solve of MeshBEMSolver3d.solve (mesh_BEM_solver_3d.py:197-373).

    For each batch entry with only Dirichlet vertices: V t = 1/2 M g + K^T g,
    returning (Neumann [B, V, 1], Dirichlet [B, V, 1]).  Neumann or mixed data
    needs the hypersingular operator, which the reference never implemented in
    3-D (its integrator raises, integrations_3d.py:440-442); the same cases
    raise NotImplementedError here."""
    if solver.BEM_type != int(BEMType.LAPLACE):
        raise NotImplementedError("solve supports BEM_type == LAPLACE")
    from .types import AssemblyMode

    dt = verts.dtype
    V = verts.shape[-2]
    B = vert_bc_flags.shape[0]
    faces = faces.to(torch.int32)
    ns = getattr(solver, "mode", int(AssemblyMode.REF)) == int(AssemblyMode.NS)
    if ns:
        if solver.order_type != int(OrderType.LINEAR):
            raise NotImplementedError("the NS Laplace rule is P1 (order_type LINEAR)")
        Vu, Ku = p1_operators(verts, faces, _lib.prec_of(dt))
    else:
        gp, gw = get_gaussion_integration_points_and_weights(solver.gaussQR, dtype=torch.float64)
        tables = relation_tables(gp.numpy(), gw.numpy())
        Vu = layer_ref(verts, faces, solver.gaussQR, 1, solver.order_type, tables)
        Ku = layer_ref(verts, faces, solver.gaussQR, 2, solver.order_type, tables)
    Vt = pack(Vu, faces, V)
    Kt = pack(Ku, faces, V)
    del Vu, Ku
    if Amat_scaler is not None:
        Vt = Amat_scaler * Vt
    va = compute_vert_areas(verts, faces, keepdim=False)
    neu, dir_ = [], []
    for b in range(B):
        flags = vert_bc_flags[b].reshape(-1)
        nD = int((flags == int(BoundaryType.DIRICHLET)).sum())
        nN = int((flags == int(BoundaryType.NEUMANN)).sum())
        if nN > 0:
            raise NotImplementedError(
                "Neumann / mixed boundary data need the hypersingular layer, absent in 3-D")
        if nD != V:
            raise ValueError(f"batch {b}: every vertex flag must be DIRICHLET or NEUMANN")
        g = vert_bc_Dirichlet[b].reshape(-1).to(dt)
        if ns:
            rhs = dirichlet_rhs(Kt, g, torch.zeros_like(va))  # K g
            rhs = mass_apply(verts, faces, g, 0.5, 1.0, rhs)  # + 1/2 M g
        else:
            rhs = dirichlet_rhs(Kt, g, va)
        t = _real_solve(Vt, rhs)
        neu.append(t.reshape(V, 1))
        dir_.append(vert_bc_Dirichlet[b].reshape(V, 1))
    return torch.stack(neu, 0), torch.stack(dir_, 0)


# ---------------------------------------------------------------- P0 (NS rule)
def p0_operators(verts: torch.Tensor, faces: torch.Tensor, prec: int, want_v: bool = True,
                 want_k: bool = True, near=None, out_v=None, out_k=None):
    """P0 Galerkin single layer V [F, F] and adjoint double layer K' [F, F]."""
    _check_vf(verts, faces)
    F = faces.shape[0]
    for o in (out_v, out_k):
        if o is not None and (o.shape != (F, F) or not o.is_contiguous()):
            raise ValueError(f"output buffers must be contiguous [F, F] (F = {F})")
    lib = _lib.load()
    dev = verts.device
    rdt = _rdt(prec)
    vt = verts.to(rdt).contiguous()
    ft = faces.to(torch.int32).contiguous()
    if near is None:
        near = near_pairs_gpu(vt, ft)
    rule = rule_tensor(dev)
    Vm = (out_v if out_v is not None else torch.empty((F, F), dtype=rdt, device=dev)) if want_v else None
    Km = (out_k if out_k is not None else torch.empty((F, F), dtype=rdt, device=dev)) if want_k else None
    wsb = lib.mba_laplace_p0_workspace_bytes(prec, F)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    st = lib.mba_laplace_p0_assemble(prec, _lib.ptr(vt), _lib.ptr(ft), verts.shape[0], F, _lib.ptr(rule),
                                     rule.shape[0], _lib.ptr(near[0]), _lib.ptr(near[1]),
                                     _lib.ptr(Vm) if Vm is not None else None,
                                     _lib.ptr(Km) if Km is not None else None, F, _lib.ptr(ws), wsb,
                                     _lib.stream())
    _lib.check(st, "mba_laplace_p0_assemble")
    return Vm, Km


# ---------------------------------------------------------------- P1 (NS rule)
def p1_operators(verts: torch.Tensor, faces: torch.Tensor, prec: int, near=None, want_v: bool = True,
                 want_k: bool = True):
    """Unpacked P1 Galerkin single and double layer [F, F, 3, 3] (layout of
    ``layer_ref``; either may be skipped)."""
    _check_vf(verts, faces)
    F = faces.shape[0]
    lib = _lib.load()
    dev = verts.device
    rdt = _rdt(prec)
    vt = verts.to(rdt).contiguous()
    ft = faces.to(torch.int32).contiguous()
    if near is None:
        near = near_pairs_gpu(vt, ft)
    rule = rule_tensor(dev)
    Vu = torch.empty((F, F, 3, 3), dtype=rdt, device=dev) if want_v else None
    Ku = torch.empty((F, F, 3, 3), dtype=rdt, device=dev) if want_k else None
    wsb = lib.mba_laplace_p0_workspace_bytes(prec, F)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    st = lib.mba_laplace_p1_assemble(prec, _lib.ptr(vt), _lib.ptr(ft), verts.shape[0], F, _lib.ptr(rule),
                                     rule.shape[0], _lib.ptr(near[0]), _lib.ptr(near[1]), _lib.ptr(Vu),
                                     _lib.ptr(Ku), _lib.ptr(ws), wsb, _lib.stream())
    _lib.check(st, "mba_laplace_p1_assemble")
    return Vu, Ku


def mass_apply(verts: torch.Tensor, faces: torch.Tensor, g: torch.Tensor, alpha: float, beta: float,
               out: torch.Tensor) -> torch.Tensor:
    """out = alpha M g + beta out with the P1 mass matrix M (in place; returns out)."""
    _check_vf(verts, faces)
    V = verts.shape[0]
    if g.numel() != V or out.numel() != V or not out.is_contiguous() or out.dtype != verts.dtype:
        raise ValueError(f"g and out must hold V = {V} values of the vertices' dtype")
    lib = _lib.load()
    vptr, vidx = vertex_incidence(faces.detach().cpu().numpy(), V)
    dev = verts.device
    st = lib.mba_laplace_mass_apply(_lib.prec_of(verts.dtype), _lib.ptr(verts.contiguous()),
                                    _lib.ptr(faces.to(torch.int32).contiguous()),
                                    _lib.ptr(torch.from_numpy(vptr).to(dev)),
                                    _lib.ptr(torch.from_numpy(vidx).to(dev)), V,
                                    _lib.ptr(g.reshape(-1).to(verts.dtype).contiguous()), float(alpha),
                                    float(beta), _lib.ptr(out), _lib.stream())
    _lib.check(st, "mba_laplace_mass_apply")
    return out
