"""Field radiation from a MoM solution: the scattered magnetic field of the
vertex currents on a set of points (``mba_radiate_h``).

The reference computes it in its demo as one dense torch expression over all
(point, vertex) pairs (``demo/demo_Helmholtz_MOM_3d.py:246-257``):

    H_scattered = -(vert_areas * cross(Js, -gradG_y(k, p1=pm, p2=verts))).sum(-2)

with ``gradG_y`` of ``src/utils/functions.py:844-868`` (wavenumber -k, so
e^{-jkR}; pairs closer than 1e-6 masked), i.e.
H(p) = sum_v A_v J_v x e^{-jkR} (1 + jkR) (p - v) / (4 pi R^3).
``radiate_H`` takes the same tensors and returns the same [Nm, 3] complex
field; the pair sum runs in one HIP kernel (LDS-staged vertex tiles, no
[Nm, V, 3] temporaries).
"""

import torch

from . import _lib


def radiate_H(wavenumber: float, points: torch.Tensor, verts: torch.Tensor, vert_Js: torch.Tensor,
              vert_areas: torch.Tensor) -> torch.Tensor:
    """Scattered H at ``points`` ([Nm, 3] or [Nm, 1, 3], as ``pm`` in the demo)
    from vertex currents ``vert_Js`` ([V, 3] or [1, V, 3] complex) on ``verts``
    ([V, 3]) with lumped vertex areas ``vert_areas`` ([V] or [V, 1]).
    Precision follows ``vert_Js`` (complex64 -> FP32, complex128 -> FP64).
    Returns complex [Nm, 3]."""
    lib = _lib.load()
    cdt = vert_Js.dtype
    if cdt not in (torch.complex64, torch.complex128):
        raise ValueError("radiate_H: vert_Js must be complex64 or complex128")
    rdt = torch.float32 if cdt == torch.complex64 else torch.float64
    pts = points.reshape(-1, 3).to(rdt).contiguous()
    V = verts.reshape(-1, 3).to(rdt).contiguous()
    J = vert_Js.reshape(-1, 3).contiguous()
    A = vert_areas.reshape(-1).to(rdt).contiguous()
    if J.shape[0] != V.shape[0] or A.shape[0] != V.shape[0]:
        raise ValueError("radiate_H: verts, vert_Js and vert_areas must have one row per vertex")
    _lib.require_device(pts, V, J, A)
    n_pts, n_verts = pts.shape[0], V.shape[0]
    out = torch.empty((n_pts, 3), dtype=cdt, device=pts.device)
    if n_pts == 0:
        return out
    if n_verts == 0:
        return out.zero_()
    prec = _lib.prec_of(cdt)
    wsb = lib.mba_radiate_h_workspace_bytes(prec, n_pts, n_verts)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=pts.device)
    st = lib.mba_radiate_h(prec, _lib.ptr(pts), n_pts, _lib.ptr(V), _lib.ptr(J), _lib.ptr(A), n_verts,
                           float(wavenumber), _lib.ptr(out), _lib.ptr(ws), wsb, _lib.stream())
    _lib.check(st, "mba_radiate_h")
    return out


def xz_grid(res: int, scale: float = 2.0, dtype=torch.float64, device="cuda") -> torch.Tensor:
    """The sampling plane of the reference demo for a square res x res grid
    (``demo_Helmholtz_MOM_3d.py:236-244``): integer meshgrid points (x = column,
    z = row, x fastest), y = 0, centred and scaled by ``mesh_normalize`` to
    [-scale, scale]^2.  Returns [res * res, 3]."""
    t = torch.linspace(-scale, scale, res, dtype=dtype, device=device)
    gz, gx = torch.meshgrid(t, t, indexing="ij")
    return torch.stack([gx.reshape(-1), torch.zeros_like(gx).reshape(-1), gz.reshape(-1)], dim=1)
