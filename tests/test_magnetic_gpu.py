"""Config-5 path on the GPU: near-pair lists, the magnetic surface-charge solve
(P0 Laplace operators in FP32 with FP64 near terms, GMRES), against the CPU FP64
oracle and the closed-form ellipsoid solution."""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd.geometry import geodesic_sphere, icosphere, noisy_ellipsoid, star_mesh
from mesh_bem_amd.magnetic import MagneticSolver, droplet_steps
from mesh_bem_amd.mom import near_pairs, near_pairs_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_near_pairs_gpu_equal_host(dev, dtype):
    for V, F in (icosphere(3), star_mesh(5, 7)[:2]):
        v = torch.tensor(V, dtype=dtype)
        p_h, i_h = near_pairs(v.double().numpy(), F)
        p_d, i_d = near_pairs_gpu(v.to(dev), torch.tensor(F, dtype=torch.int32, device=dev))
        assert np.array_equal(p_d.cpu().numpy(), p_h)
        assert np.array_equal(i_d[: p_h[-1]].cpu().numpy(), i_h)


def test_magnetic_matches_oracle(dev):
    V, F = geodesic_sphere(6)
    V = V * np.array([1.0, 1.0, 1.4])
    chi = 2.5
    v = torch.tensor(V, dtype=torch.float64, device=dev)
    f = torch.tensor(F, dtype=torch.int32, device=dev)
    ms = MagneticSolver(chi, tol=1e-9)
    out = ms.solve(v, f)
    assert out.info.converged
    ptr, idx = O.near_pairs_bruteforce(V, F)
    Vo, Ko = O.laplace_p0(V.astype(np.float32).astype(np.float64), F, ptr, idx)
    A, n = O.face_areas_normals(V, F)
    sig = np.linalg.solve((1 / chi + 0.5) * np.diag(A) + Ko, A * n[:, 2])
    e = np.linalg.norm(out.sigma.cpu().numpy() - sig) / np.linalg.norm(sig)
    phi = (Vo @ sig) / A
    ep = np.linalg.norm(out.phi.cpu().numpy() - phi) / np.linalg.norm(phi)
    print(f"  sigma relerr {e:.2e}, phi relerr {ep:.2e}, iterations {out.info.iterations}")
    assert e < 2e-5 and ep < 2e-5


def test_magnetic_ellipsoid_closed_form(dev):
    c = 1.5
    V, F = geodesic_sphere(14)
    V = V * np.array([1.0, 1.0, c])
    chi = 3.0
    out = MagneticSolver(chi).solve(torch.tensor(V, device=dev), torch.tensor(F, dtype=torch.int32, device=dev))
    _, _, Nc = O.demag_factors(1.0, 1.0, c)
    A, n = O.face_areas_normals(V, F)
    ref = chi / (1 + Nc * chi) * n[:, 2]
    e = np.linalg.norm(out.sigma.cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"  ellipsoid c={c} chi={chi} F={len(F)}: relerr vs closed form {e:.3e}")
    assert out.info.converged and e < 0.02


def test_warm_start_cuts_iterations(dev):
    ms = MagneticSolver(4.0, tol=1e-6, with_potential=False)
    prev = None
    its = []
    for s, V, F, c in droplet_steps(3, 10, 3):
        v = torch.tensor(V, device=dev)
        f = torch.tensor(F, dtype=torch.int32, device=dev)
        cold = ms.solve(v, f)
        if prev is not None:
            warm = ms.solve(v, f, x0=prev)
            its.append((cold.info.iterations, warm.info.iterations))
            d = np.linalg.norm(warm.sigma.cpu().numpy() - cold.sigma.cpu().numpy()) / np.linalg.norm(
                cold.sigma.cpu().numpy())
            assert warm.info.converged and d < 1e-5
        prev = cold.sigma
    print(f"  (cold, warm) iterations: {its}")
    assert all(w < c for c, w in its)


def test_droplet_step_with_magnetic_pressure(dev):
    """One step of the ferrofluid loop on a resting sphere in a uniform field
    (no gravity, no surface tension): the magnetic traction is largest at the
    poles, so the pressure solve accelerates the poles outwards and the
    equator inwards (the droplet starts to elongate along H0).  Later steps
    warm-start the magnetic GMRES and converge."""
    from mesh_bem_amd.magnetic import FerrofluidDroplet
    from mesh_bem_amd.types import AssemblyMode

    V, F = geodesic_sphere(8)
    v = torch.tensor(V, dtype=torch.float64, device=dev)
    f = torch.tensor(F, dtype=torch.int32, device=dev)
    vel = torch.zeros((1, len(V), 3), dtype=torch.float64, device=dev)
    drop = FerrofluidDroplet(chi=3.0, gaussQR=3, device=dev, dtype=torch.float64, bem_mode=int(AssemblyMode.NS))
    v1, f1, vel1, rec = drop.step(v, f, vel, dt=1e-2)
    n = torch.tensor(V / np.linalg.norm(V, axis=1, keepdims=True), device=dev)
    un = (vel1[0] * n).sum(-1)
    poles, equator = n[:, 2].abs() > 0.9, n[:, 2].abs() < 0.1
    print(f"  step 1: {rec}; outward velocity poles {float(un[poles].mean()):.3e} "
          f"equator {float(un[equator].mean()):.3e}")
    assert torch.isfinite(vel1).all() and rec["gmres_converged"] and not rec["warm_start"]
    assert float(un[poles].min()) > 0 and float(un[equator].max()) < 0
    its = [rec["gmres_iterations"]]
    for _ in range(2):
        v1, f1, vel1, rec = drop.step(v1, f1, vel1, dt=1e-2)
        assert rec["warm_start"] and rec["gmres_converged"] and torch.isfinite(vel1).all()
        its.append(rec["gmres_iterations"])
    print(f"  magnetic GMRES iterations per step {its}")
    assert its[2] < its[0]
