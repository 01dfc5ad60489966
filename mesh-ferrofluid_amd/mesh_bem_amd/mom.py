"""RWG MoM assembly driver (host side).

The reference assembles face-edge blocks on the GPU, then scatters them into
RWG space with ``scatter_add_`` (src/mesh_BEM/mesh_BEM_solver_3d.py:128-164).
Here the HIP library writes face-edge blocks for a window of source faces and
a fixed-order gather kernel accumulates them into the row-major RWG matrix,
so the result is deterministic and the face-edge buffer is bounded by the
window size instead of being 9 F^2 entries.

Topology (integer work) runs on the host in numpy:
* RWG face-edge pairs ``rwg_fe`` [E,2] from ``face_edge_indices``;
* near-pair lists: test face a is "near" source face b when
  |o_a - o_b| < NEAR_ETA * (rho_a + rho_b) (o = centroid, rho = circumradius
  about o); every pair sharing a vertex satisfies it;
* vertex incidence lists for the deterministic vertex sum.
"""

from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib
from .geometry import check_mesh, check_orientation, get_edge_unique, rwg_face_edges
from .quadrature import radon7

NEAR_ETA = 2.0


def near_pairs(verts: np.ndarray, faces: np.ndarray, eta: float = NEAR_ETA):
    """CSR over source faces b of the near test faces a (ascending)."""
    from scipy.spatial import cKDTree

    v = np.asarray(verts, dtype=np.float64)
    f = np.asarray(faces, dtype=np.int64)
    P = v[f]
    o = P.mean(axis=1)
    rho = np.linalg.norm(P - o[:, None, :], axis=2).max(axis=1)
    tree = cKDTree(o)
    rmax = rho.max()
    cand = tree.query_ball_point(o, r=eta * (rho + rmax) * (1 + 1e-12))
    ptr = np.zeros(len(f) + 1, dtype=np.int64)
    idx = []
    for b, lst in enumerate(cand):
        a = np.asarray(sorted(lst), dtype=np.int64)
        d = np.linalg.norm(o[a] - o[b], axis=1)
        a = a[d < eta * (rho[a] + rho[b])]
        idx.append(a)
        ptr[b + 1] = ptr[b] + a.size
    return ptr, np.concatenate(idx).astype(np.int32)


def near_pairs_gpu(verts: torch.Tensor, faces: torch.Tensor, eta: float = NEAR_ETA):
    """Near-pair CSR built on the GPU (same criterion and order as ``near_pairs``)."""
    lib = _lib.load()
    vt = verts.contiguous()
    ft = faces.to(torch.int32).contiguous()
    F = ft.shape[0]
    dev = vt.device
    wsb = lib.mba_near_workspace_bytes(F)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    ptr = torch.empty(F + 1, dtype=torch.int64, device=dev)
    st = lib.mba_near_pairs_count(_lib.prec_of(vt.dtype), _lib.ptr(vt), _lib.ptr(ft), vt.shape[0], F,
                                  float(eta), _lib.ptr(ptr), _lib.ptr(ws), wsb, _lib.stream())
    _lib.check(st, "mba_near_pairs_count")
    nnz = int(ptr[F].item())
    idx = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    st = lib.mba_near_pairs_fill(F, float(eta), _lib.ptr(ptr), _lib.ptr(idx), _lib.ptr(ws), wsb, _lib.stream())
    _lib.check(st, "mba_near_pairs_fill")
    return ptr, idx


def vertex_incidence(faces: np.ndarray, n_verts: int):
    """CSR vertex -> face-vertex slots (3f + t), ascending slot order."""
    f = np.asarray(faces, dtype=np.int64).reshape(-1)
    order = np.argsort(f, kind="stable")
    ptr = np.zeros(n_verts + 1, dtype=np.int64)
    np.add.at(ptr, f + 1, 1)
    ptr = np.cumsum(ptr)
    return ptr, order.astype(np.int32)


def rule_tensor(device, rule=None) -> torch.Tensor:
    lam, w = radon7() if rule is None else rule
    tab = np.concatenate([np.asarray(lam, dtype=np.float64), np.asarray(w, dtype=np.float64)[:, None]], 1)
    return torch.from_numpy(np.ascontiguousarray(tab)).to(device)


@dataclass
class MomMesh:
    """Device-resident mesh with the tables the MoM kernels use."""

    verts: torch.Tensor  # [V,3] float/double, device
    faces: torch.Tensor  # [F,3] int32, device
    face_edge_indices: torch.Tensor  # [3F] int32, device
    n_edges: int
    rwg_fe: torch.Tensor  # [E,2] int32, device
    near_ptr: torch.Tensor  # [F+1] int64
    near_idx: torch.Tensor  # [nnz] int32
    vinc_ptr: torch.Tensor
    vinc_idx: torch.Tensor
    extra: dict = field(default_factory=dict)

    @property
    def n_faces(self):
        return self.faces.shape[0]

    @property
    def n_verts(self):
        return self.verts.shape[0]

    @staticmethod
    def build(verts, faces, face_edge_indices=None, n_edges=None, device=None, dtype=None,
              eta: float = NEAR_ETA, near: bool = True) -> "MomMesh":
        device = device or (verts.device if isinstance(verts, torch.Tensor) else torch.device("cuda"))
        vt = torch.as_tensor(verts)
        if dtype is not None:
            vt = vt.to(dtype)
        ft = torch.as_tensor(faces).to(torch.int32)
        vnp = vt.detach().cpu().numpy().astype(np.float64)
        fnp = ft.detach().cpu().numpy().astype(np.int64)
        if vnp.ndim != 2 or vnp.shape[1] != 3 or fnp.ndim != 2 or fnp.shape[1] != 3 or len(fnp) == 0:
            raise ValueError("MomMesh.build needs verts [V, 3] and faces [F, 3], F > 0")
        if fnp.min() < 0 or fnp.max() >= len(vnp):
            raise ValueError("face vertex index out of range")
        if face_edge_indices is None:
            edges, fe, _ = get_edge_unique(vt.cpu(), ft.cpu())
            n_edges = edges.shape[0]
            face_edge_indices = fe
        fe_np = torch.as_tensor(face_edge_indices).detach().cpu().numpy().astype(np.int64).reshape(-1)
        if n_edges is None:
            n_edges = int(fe_np.max()) + 1
        rfe = rwg_face_edges(fe_np, n_edges)
        check_mesh(vnp, fnp)
        check_orientation(fnp, rfe)
        vdev = vt.to(device).contiguous()
        fdev = ft.to(device).contiguous()
        if not near:
            nptr = torch.zeros(len(fnp) + 1, dtype=torch.int64, device=device)
            nidx = torch.zeros(1, dtype=torch.int32, device=device)
        elif torch.device(device).type == "cuda":
            nptr, nidx = near_pairs_gpu(vdev, fdev, eta)
        else:
            p, i = near_pairs(vnp, fnp, eta)
            nptr, nidx = torch.from_numpy(p).to(device), torch.from_numpy(i).to(device)
        vptr, vidx = vertex_incidence(fnp, vnp.shape[0])
        return MomMesh(
            verts=vdev,
            faces=fdev,
            face_edge_indices=torch.from_numpy(fe_np.astype(np.int32)).to(device),
            n_edges=int(n_edges),
            rwg_fe=torch.from_numpy(rfe).to(device),
            near_ptr=nptr,
            near_idx=nidx,
            vinc_ptr=torch.from_numpy(vptr).to(device),
            vinc_idx=torch.from_numpy(vidx).to(device),
        )


def _cdt(prec: int):
    return torch.complex64 if prec == _lib.MBA_F32 else torch.complex128


def _rdt(prec: int):
    return torch.float32 if prec == _lib.MBA_F32 else torch.float64


def zmat_faceedge_ns(mesh: MomMesh, k: float, prec: int, b0: int = 0, nb: Optional[int] = None,
                     out: Optional[torch.Tensor] = None, rule: Optional[torch.Tensor] = None):
    """Face-edge EFIE blocks for source faces [b0, b0+nb): [3F, 3nb] complex."""
    lib = _lib.load()
    F = mesh.n_faces
    nb = F - b0 if nb is None else nb
    dev = mesh.verts.device
    if rule is None:
        rule = rule_tensor(dev)
    verts = mesh.verts.to(_rdt(prec)).contiguous()
    if out is None:
        out = torch.empty((3 * F, 3 * nb), dtype=_cdt(prec), device=dev)
    ws_bytes = lib.mba_mom_workspace_bytes(prec, F)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    st = lib.mba_mom_zmat_faceedge_ns(
        prec, _lib.ptr(verts), _lib.ptr(mesh.faces), mesh.n_verts, F, _lib.ptr(rule), rule.shape[0],
        float(k), _lib.ptr(mesh.near_ptr), _lib.ptr(mesh.near_idx), b0, nb, _lib.ptr(out),
        out.stride(0), _lib.ptr(ws), ws_bytes, _lib.stream())
    _lib.check(st, "mba_mom_zmat_faceedge_ns")
    return out


def assemble_ns(mesh: MomMesh, k: float, prec: int, window_faces: Optional[int] = None,
                out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """RWG-space EFIE matrix Z [E, E] (row = test RWG), complex64/complex128.

    Source faces are processed in windows of ``window_faces`` (default: the
    whole mesh when the face-edge buffer stays under 64 GiB, so every entry of
    Z is written once and in order; a column whose two face-edges fall in
    different windows is stored, then added to)."""
    lib = _lib.load()
    F, E = mesh.n_faces, mesh.n_edges
    dev = mesh.verts.device
    cdt = _cdt(prec)
    esz = 8 if prec == _lib.MBA_F32 else 16
    if window_faces is None:
        window_faces = max(64, min(F, int((64 << 30) // (9 * F * esz))))
    window_faces = min(window_faces, F)
    if out is None:
        out = torch.empty((E, E), dtype=cdt, device=dev)
    zf = torch.empty((3 * F, 3 * window_faces), dtype=cdt, device=dev)
    rule = rule_tensor(dev)
    for b0 in range(0, F, window_faces):
        nb = min(window_faces, F - b0)
        zfw = zf[:, : 3 * nb] if nb < window_faces else zf
        zmat_faceedge_ns(mesh, k, prec, b0, nb, out=zfw, rule=rule)
        st = lib.mba_mom_gather_rwg(prec, _lib.ptr(zf), zf.stride(0), 3 * b0, 3 * nb,
                                    _lib.ptr(mesh.rwg_fe), E, _lib.ptr(out), out.stride(0),
                                    _lib.stream())
        _lib.check(st, "mba_mom_gather_rwg")
    return out


def rhs_ns(mesh: MomMesh, k: float, prec: int, dirs, pols) -> torch.Tensor:
    """Plane-wave RHS V [E, ndir] (complex)."""
    dev = mesh.verts.device
    dirs = torch.as_tensor(np.atleast_2d(np.asarray(dirs, dtype=np.float64))).to(dev).contiguous()
    pols = torch.as_tensor(np.atleast_2d(np.asarray(pols, dtype=np.float64))).to(dev).contiguous()
    return _rhs_ns_dev(mesh, k, prec, dirs, pols, rule_tensor(dev))


def rhs_ns_many(meshes, ks, prec: int, dirs_list, pols_list) -> list:
    """rhs_ns for many meshes: the directions and polarisations of all of
    them and the quadrature rule go to the device in one copy each."""
    if not meshes:
        return []
    dev = meshes[0].verts.device
    ds = [np.atleast_2d(np.asarray(d, dtype=np.float64)) for d in dirs_list]
    ps = [np.atleast_2d(np.asarray(p, dtype=np.float64)) for p in pols_list]
    if any(d.shape != p.shape or d.shape[1] != 3 for d, p in zip(ds, ps)):
        raise ValueError("rhs_ns_many: dirs and pols must be [ndir, 3] per mesh")
    D = torch.from_numpy(np.ascontiguousarray(np.concatenate(ds))).to(dev)
    P = torch.from_numpy(np.ascontiguousarray(np.concatenate(ps))).to(dev)
    rule = rule_tensor(dev)
    out, off = [], 0
    for m, k, d in zip(meshes, ks, ds):
        nd = d.shape[0]
        out.append(_rhs_ns_dev(m, k, prec, D[off:off + nd], P[off:off + nd], rule))
        off += nd
    return out


def _rhs_ns_dev(mesh: MomMesh, k: float, prec: int, dirs: torch.Tensor, pols: torch.Tensor,
                rule: torch.Tensor) -> torch.Tensor:
    """rhs_ns with device-resident contiguous dirs / pols [ndir, 3] (FP64) and rule."""
    lib = _lib.load()
    dev = mesh.verts.device
    nd = dirs.shape[0]
    F, E = mesh.n_faces, mesh.n_edges
    verts = mesh.verts.to(_rdt(prec)).contiguous()
    vf = torch.empty((3 * F, nd), dtype=_cdt(prec), device=dev)
    st = lib.mba_mom_rhs_faceedge_ns(prec, _lib.ptr(verts), _lib.ptr(mesh.faces), mesh.n_verts, F,
                                     _lib.ptr(rule), rule.shape[0], float(k), _lib.ptr(dirs),
                                     _lib.ptr(pols), nd, _lib.ptr(vf), _lib.stream())
    _lib.check(st, "mba_mom_rhs_faceedge_ns")
    v = torch.empty((E, nd), dtype=_cdt(prec), device=dev)
    st = lib.mba_mom_gather_rhs(prec, _lib.ptr(vf), _lib.ptr(mesh.rwg_fe), E, nd, _lib.ptr(v),
                                _lib.stream())
    _lib.check(st, "mba_mom_gather_rhs")
    return v


def currents_to_vertices(mesh: MomMesh, icoef: torch.Tensor, face_areas: torch.Tensor = None):
    """RWG coefficients [B, E] -> face-vertex currents [B, 3F, 3] and vertex sums [B, V, 3]."""
    from .geometry import compute_face_areas

    lib = _lib.load()
    dt = mesh.verts.dtype
    prec = _lib.prec_of(dt)
    B = icoef.shape[0]
    F, V = mesh.n_faces, mesh.n_verts
    if face_areas is None:
        face_areas = compute_face_areas(mesh.verts, mesh.faces, keepdim=False)
    fa = face_areas.reshape(-1).to(dt).contiguous()
    ic = icoef.reshape(B, -1).to(_cdt(prec)).contiguous()
    fvj = torch.empty((B, 3 * F, 3), dtype=_cdt(prec), device=mesh.verts.device)
    st = lib.mba_mom_interp_faceedge(prec, _lib.ptr(mesh.verts), _lib.ptr(mesh.faces),
                                     _lib.ptr(mesh.face_edge_indices), _lib.ptr(fa), _lib.ptr(ic),
                                     B, ic.shape[1], V, F, _lib.ptr(fvj), _lib.stream())
    _lib.check(st, "mba_mom_interp_faceedge")
    vj = torch.empty((B, V, 3), dtype=_cdt(prec), device=mesh.verts.device)
    st = lib.mba_mom_vertex_sum(prec, _lib.ptr(fvj), _lib.ptr(mesh.vinc_ptr),
                                _lib.ptr(mesh.vinc_idx), B, V, F, _lib.ptr(vj), _lib.stream())
    _lib.check(st, "mba_mom_vertex_sum")
    return fvj, vj
