"""Magnetic scalar potential of a linear magnetisable surface (Laplace BEM, P0).

The reference stores a susceptibility but has no magnetic solver; its time
step only exposes a ``P_mag`` hook (src/mesh_ferrofluid/mesh_pressure.py:109-110,
src/simulation/simulation.py:110).  This module supplies it on the P0 Laplace
operators.

A body of susceptibility chi in a uniform applied field H0 carries the surface
charge sigma = M . n; with H = H0 - grad(S sigma) (S the single layer) the
interior normal field gives, in Galerkin P0 form with the lumped mass
M = diag(area),
    [ (1/chi + 1/2) M + K' ] sigma = M (H0 . n),
K' the adjoint double layer.  The potential on the surface is phi = S sigma
(face means (V sigma)_i / area_i).  For an ellipsoid sigma = chi / (1 + N_z chi)
n_z H0 exactly (N_z its demagnetisation factor), which pins the formulation.

Per step: FP32 assembly of V and K' (7-point rule, FP64 analytic near terms,
near-pair lists rebuilt on the GPU), the diagonal shift in place, and GMRES with
an FP32 matrix and FP64 iterate, warm-started from the previous step's charge.
"""

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .geometry import compute_face_areas, compute_face_normals, noisy_ellipsoid
from .gmres import GmresInfo, gmres, matvec
from .laplace import p0_operators


@dataclass
class MagneticStep:
    sigma: torch.Tensor  # [F] float64 surface charge
    phi: Optional[torch.Tensor]  # [F] float64 face-mean potential (None if V not assembled)
    info: GmresInfo
    assembly_s: float = 0.0
    solve_s: float = 0.0


class MagneticSolver:
    """Surface charge of a linear magnetisable closed surface in a uniform field."""

    def __init__(self, chi: float, H0: Sequence[float] = (0.0, 0.0, 1.0), restart: int = 50,
                 tol: float = 1e-6, with_potential: bool = True):
        if not chi > 0:
            raise ValueError("susceptibility must be positive")
        self.chi = float(chi)
        self.H0 = np.asarray(H0, dtype=np.float64)
        self.restart = restart
        self.tol = tol
        self.with_potential = with_potential
        self._buf = {}

    def _buffers(self, F, dev):
        key = (F, str(dev))
        if key not in self._buf:
            self._buf = {key: (torch.empty((F, F), dtype=torch.float32, device=dev),
                               torch.empty((F, F), dtype=torch.float32, device=dev)
                               if self.with_potential else None)}
        return self._buf[key]

    def system(self, verts: torch.Tensor, faces: torch.Tensor):
        """(A [F,F] float32 = lam M + K', b [F] float64, V [F,F] float32 or None, areas)."""
        lib = _lib.load()
        F = faces.shape[0]
        dev = verts.device
        A, Vb = self._buffers(F, dev)
        vt = verts.to(torch.float32).contiguous()
        ft = faces.to(torch.int32).contiguous()
        Vm, Km = p0_operators(vt, ft, _lib.MBA_F32, want_v=Vb is not None, want_k=True, out_v=Vb, out_k=A)
        v64 = verts.to(torch.float64)
        area = compute_face_areas(v64, ft, keepdim=False).contiguous()
        nrm = compute_face_normals(v64, ft)
        lam = 1.0 / self.chi + 0.5
        a32 = area.to(torch.float32).contiguous()
        st = lib.mba_add_diag(_lib.MBA_F32, _lib.ptr(Km), F, Km.stride(0), lam, _lib.ptr(a32), _lib.stream())
        _lib.check(st, "mba_add_diag")
        b = area * (nrm @ torch.as_tensor(self.H0, dtype=torch.float64, device=dev))
        return Km, b.contiguous(), Vm, area

    def solve(self, verts: torch.Tensor, faces: torch.Tensor, x0: Optional[torch.Tensor] = None,
              maxiter: int = 1000) -> MagneticStep:
        import time

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        A, b, Vm, area = self.system(verts, faces)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # a non-finite warm start restarts from zero; an unconverged GMRES is
        # finished by LU (both recorded in info.fallback)
        sigma, info = gmres(A, b, x0=x0, restart=self.restart, tol=self.tol, maxiter=maxiter, fallback="lu")
        phi = matvec(Vm, sigma) / area if Vm is not None else None
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return MagneticStep(sigma, phi, info, t1 - t0, t2 - t1)


def droplet_steps(seed: int, freq: int, steps: int, noise: float = 0.02, step_noise: float = 0.002):
    """Vertex sets of a deforming droplet: a noisy ellipsoid (axes 1:1:c,
    c ~ U(1,2), radial noise N(0, noise^2)) followed by ``steps`` radial
    random-walk perturbations N(0, step_noise^2) (numpy PCG64, seeded per step).
    Yields (step, verts [V,3] float64, faces [F,3], c)."""
    V0, F0, c = noisy_ellipsoid(seed, freq, noise)
    nrm = V0 / np.linalg.norm(V0, axis=1, keepdims=True)
    V = V0.copy()
    yield 0, V.copy(), F0, c
    for s in range(1, steps):
        rng = np.random.Generator(np.random.PCG64(seed * 100003 + s))
        V = V + nrm * rng.normal(0.0, step_noise, size=(V.shape[0], 1))
        yield s, V.copy(), F0, c


# ---------------------------------------------------------------- magnetic pressure
def magnetic_pressure(verts: torch.Tensor, faces: torch.Tensor, sigma: torch.Tensor, phi: torch.Tensor,
                      chi: float, H0: Sequence[float] = (0.0, 0.0, 1.0), mu0: float = 1.0) -> torch.Tensor:
    """Magnetic normal traction at the vertices, the ``P_mag`` of the
    reference's pressure step (src/mesh_ferrofluid/mesh_pressure.py:108-110,
    subtracted from the Dirichlet pressure there).

    For a linear fluid (M = chi H) the traction on the interface is
        P_mag = mu0 / 2 [ (M . n)^2 + chi |H_in|^2 ]
    (the jump of the Maxwell stress plus the magnetic fluid pressure
    mu0 int_0^H M dH of a linear material), with H_in the interior field:
    normal part H_in . n = sigma / chi (sigma = M . n the surface charge),
    tangential part (I - n n^T)(H0 - grad phi) (phi = S sigma, continuous
    across the surface).  sigma and phi are P0 face values (MagneticSolver);
    phi is averaged to the vertices (area weights), differentiated on each
    face, and the face values of P_mag are averaged back to the vertices.
    For an ellipsoid, H_in is uniform, H0 / (1 + N_z chi), which pins the
    result: P_mag = mu0 / 2 chi |H_in|^2 (chi n_z^2 + 1)."""
    v = verts.to(torch.float64)
    f = faces.to(torch.int64)
    V, F = v.shape[0], f.shape[0]
    P = v[f]  # [F, 3, 3]
    cr = torch.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 1], dim=-1)
    a2 = cr.norm(dim=-1)
    n = cr / a2[:, None]
    area = 0.5 * a2
    sig = sigma.to(torch.float64).reshape(F)
    ph = phi.to(torch.float64).reshape(F)
    # vertex potential: area-weighted mean of the incident faces
    wsum = torch.zeros(V, dtype=torch.float64, device=v.device).index_add_(0, f.reshape(-1), area.repeat_interleave(3))
    phv = torch.zeros(V, dtype=torch.float64, device=v.device).index_add_(
        0, f.reshape(-1), (area * ph).repeat_interleave(3)) / wsum
    # gradient of the linear interpolant on each face: sum_i phi_i n x (P_i+2 - P_i+1) / (2A)
    grad = torch.zeros((F, 3), dtype=torch.float64, device=v.device)
    for i in range(3):
        e = P[:, (i + 2) % 3] - P[:, (i + 1) % 3]
        grad += phv[f[:, i]][:, None] * torch.cross(n, e, dim=-1) / a2[:, None]
    h0 = torch.as_tensor(np.asarray(H0, dtype=np.float64), device=v.device)
    Ht = h0[None, :] - grad
    Ht = Ht - (Ht * n).sum(-1, keepdim=True) * n
    Hn = sig / chi
    pf = 0.5 * mu0 * (sig * sig + chi * ((Ht * Ht).sum(-1) + Hn * Hn))
    pv = torch.zeros(V, dtype=torch.float64, device=v.device).index_add_(0, f.reshape(-1),
                                                                          (area * pf).repeat_interleave(3))
    return (pv / wsum).to(verts.dtype)


class MeshAdvection:
    """Drop-in for src/mesh_ferrofluid/mesh_advection.py:4-22: verts + vel * dt."""

    def solve(self, verts: torch.Tensor, dt: float, vel: torch.Tensor) -> torch.Tensor:
        return verts + vel * dt


class FerrofluidDroplet:
    """One time step of the surface-only ferrofluid droplet loop
    (demo/demo_3d_mesh_droplet.py:134-164) with the magnetic pressure the
    reference leaves out (it passes P_mag=None): advection, optional remesh,
    Helmholtz decomposition, the magnetic surface charge (P0 Laplace BEM,
    GMRES warm-started from the previous step while the mesh keeps its
    faces), P_mag, and the pressure solve with P_mag."""

    def __init__(self, chi: float, density_fluid: float = 1.0, gravity_strength: float = 0.0,
                 surface_tension: float = 0.0, gaussQR: int = 3, order_type: int = 1,
                 H0: Sequence[float] = (0.0, 0.0, 1.0), mu0: float = 1.0, target_edge_length: Optional[float] = None,
                 device=torch.device("cuda"), dtype: torch.dtype = torch.float64, bem_mode: Optional[int] = None,
                 tol: float = 1e-6):
        from .ferrofluid import MeshHelmholtzDecomposition, MeshPressure

        self.chi = float(chi)
        self.H0 = tuple(float(x) for x in H0)
        self.mu0 = float(mu0)
        self.target_edge_length = target_edge_length
        self.advection = MeshAdvection()
        self.decomposition = MeshHelmholtzDecomposition(gaussQR, device=device, dtype=dtype)
        self.pressure = MeshPressure(density_fluid, gravity_strength, surface_tension, 0.5 * np.pi, gaussQR,
                                     order_type, device=device, dtype=dtype, bem_mode=bem_mode)
        self.magnetic = MagneticSolver(chi, H0=self.H0, tol=tol, with_potential=True)
        self._sigma = None

    def step(self, verts: torch.Tensor, faces: torch.Tensor, vel: torch.Tensor, dt: float):
        """verts [V, 3], faces [F, 3], vel [B, V, 3] -> (verts, faces, vel, record)."""
        from .remesh import remesh_and_transfer_velocity
        from .types import BoundaryType

        verts = self.advection.solve(verts, dt, vel[0])
        if self.target_edge_length is not None:
            verts, faces, vel = remesh_and_transfer_velocity(verts, faces, vel, self.target_edge_length)
            self._sigma = None
        flags = torch.full((vel.shape[0], verts.shape[0], 1), int(BoundaryType.DIRICHLET), dtype=torch.uint8,
                           device=verts.device)
        vel = self.decomposition.solve(verts, faces, vel)
        x0 = self._sigma if self._sigma is not None and self._sigma.numel() == faces.shape[0] else None
        mag = self.magnetic.solve(verts, faces, x0=x0)
        self._sigma = mag.sigma
        p_mag = magnetic_pressure(verts, faces, mag.sigma, mag.phi, self.chi, self.H0, self.mu0)
        vel = self.pressure.solve(verts, faces, dt, vel, flags, P_mag=p_mag.reshape(1, -1, 1))
        rec = {"gmres_iterations": mag.info.iterations, "gmres_converged": mag.info.converged,
               "gmres_fallback": mag.info.fallback, "warm_start": x0 is not None,
               "p_mag_max": float(p_mag.max()), "p_mag_min": float(p_mag.min())}
        return verts, faces, vel, rec
