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
