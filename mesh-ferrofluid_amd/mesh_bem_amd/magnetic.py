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
